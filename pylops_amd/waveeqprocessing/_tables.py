"""Traveltime / amplitude / angle tables, built in float64 with torch on the
operator's device, and their encoding into the tiled device layout of
libkirchx (kx_encode_table / kx_tile_table).

Table semantics follow pylops/waveeqprocessing/kirchhoff.py:
analytic traveltimes K:570-607, gradients K:642-648 (and unit spacing for
byot, K:351-360), amplitudes K:316-330, angles K:364-393, six/rix
K:289-301, apertures K:414-439.
"""
import ctypes
import math

import numpy as np
import torch

from pylops_amd import _lib

TAPER_LEN = 21


def as_f64(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64)
    return torch.as_tensor(np.asarray(a, dtype=np.float64), device=device)


def div(a, b, rounding_mode=None):
    """a / b with b a Python scalar, as an IEEE division of every element.
    (torch turns division by a scalar into multiplication by its rounded
    reciprocal, which is off by one ulp against numpy for some inputs.)"""
    bt = torch.tensor(float(b), dtype=a.dtype, device=a.device).expand_as(a)
    return torch.div(a, bt, rounding_mode=rounding_mode)


def grid_points(z, x, y, device):
    """Raveled coordinates of the image points in reference order
    (meshgrid indexing="ij", z fastest)."""
    xz = torch.as_tensor(np.asarray(z, dtype=np.float64), device=device)
    xx = torch.as_tensor(np.asarray(x, dtype=np.float64), device=device)
    if y is None:
        X, Z = torch.meshgrid(xx, xz, indexing="ij")
        return X.reshape(-1), Z.reshape(-1), None
    yy = torch.as_tensor(np.asarray(y, dtype=np.float64), device=device)
    Y, X, Z = torch.meshgrid(yy, xx, xz, indexing="ij")
    return X.reshape(-1), Z.reshape(-1), Y.reshape(-1)


def analytic(z, x, srcs, recs, vel, y, dims, dsamp, device):
    """(trav_srcs, trav_recs, dist_srcs, dist_recs, grad_srcs, grad_recs) for
    a constant velocity; all float64 torch tensors of shape (ni, n)."""
    X, Z, Y = grid_points(z, x, y, device)
    sd = 0 if y is None else 1

    def dist2(pts):
        p = as_f64(pts, device)
        dx = X[:, None] - p[0 + sd][None, :]
        dz = Z[:, None] - p[1 + sd][None, :]
        d2 = dx * dx + dz * dz
        if Y is not None:
            dy = Y[:, None] - p[0][None, :]
            d2 = d2 + dy * dy
        return d2

    ts = div(torch.sqrt(dist2(srcs)), vel)
    tr = div(torch.sqrt(dist2(recs)), vel)
    ds = ts * vel
    dr = tr * vel
    return ts, tr, ds, dr, gradient(ts, dims, dsamp), gradient(tr, dims, dsamp)


def eikonal(pts, vel, origin, dsamp, dims, device, unit=False, max_iter=20, order=2):
    """Traveltime (or, with ``unit``, distance) tables of shape (ni, n) from
    the grid cells nearest to the points ``pts`` ((x, z) or (y, x, z) rows,
    rounded as K:617-633 rounds them), by the device eikonal solver
    (``kx_eikonal``; ``order`` 2 is scikit-fmm's default).  ``vel`` is a
    scalar or an array of the grid's shape."""
    import ctypes
    ndims = len(dims)
    pts = np.asarray(pts, dtype=np.float64)
    cells = np.round((pts.T - np.asarray(origin)) / np.asarray(dsamp)).astype(np.int32)
    if np.any(cells < 0) or np.any(cells >= np.asarray(dims)):
        raise ValueError("every source and receiver must lie inside the image grid "
                         "for mode='eikonal'")
    if ndims == 2:
        (nx, nz), ny = dims, 1
        hx, hz, hy = float(dsamp[0]), float(dsamp[1]), 1.0
        src = np.stack((cells[:, 0], np.zeros(len(cells), np.int32), cells[:, 1]), axis=1)
    else:
        ny, nx, nz = dims
        hy, hx, hz = (float(v) for v in dsamp)
        src = np.stack((cells[:, 1], cells[:, 0], cells[:, 2]), axis=1)
    ni = ny * nx * nz
    if unit:
        slow = torch.ones(ni, dtype=torch.float64, device=device)
    else:
        v = as_f64(vel, device)
        v = v.expand(ni) if v.numel() == 1 else v.reshape(-1)
        if v.numel() != ni:
            raise ValueError(f"vel of {v.numel()} values for a grid of {ni} points")
        slow = 1.0 / v
    n = len(src)
    srcd = torch.as_tensor(np.ascontiguousarray(src, dtype=np.int32), device=device)
    out = torch.empty((ni, n), dtype=torch.float64, device=device)
    flag = torch.zeros(1, dtype=torch.int32, device=device)
    iters = (ctypes.c_int32 * 2)(0, 0)
    L = _lib.load()
    rc = L.kx_eikonal(ndims, nx, ny, nz, hx, hy, hz, _lib.ptr(slow.contiguous()), n,
                      _lib.ptr(srcd), _lib.ptr(out), max_iter, order, _lib.ptr(flag),
                      iters, _lib.stream_ptr(device))
    _lib.check(rc, "kx_eikonal")
    if not iters[1]:
        # an unconverged sweep leaves upper bounds, not first arrivals
        raise RuntimeError(
            f"eikonal traveltimes did not converge in {max_iter} sweep rounds per phase "
            f"({iters[0]} rounds run); pass a larger eikonal_iter")
    return out


def eikonal_tables(srcs, recs, vel, origin, dsamp, dims, device, distances, max_iter=20):
    """(trav_srcs, trav_recs, dist_srcs, dist_recs, grad_srcs, grad_recs) of
    the eikonal mode (K:609-657): first-arrival times from each source and
    receiver cell, the unit-speed distances when ``distances`` (dynamic
    amplitudes need them), and the traveltime gradients; ``max_iter`` sweep
    rounds per phase at most (RuntimeError if they do not converge)."""
    ts = eikonal(srcs, vel, origin, dsamp, dims, device, max_iter=max_iter)
    tr = eikonal(recs, vel, origin, dsamp, dims, device, max_iter=max_iter)
    ds = dr = None
    if distances:
        ds = eikonal(srcs, None, origin, dsamp, dims, device, unit=True, max_iter=max_iter)
        dr = eikonal(recs, None, origin, dsamp, dims, device, unit=True, max_iter=max_iter)
    return ts, tr, ds, dr, gradient(ts, dims, dsamp), gradient(tr, dims, dsamp)


def gradient(tab, dims, dsamp=None):
    """np.gradient(tab.reshape(*dims, n), *dsamp, axis=range(ndims)) with
    edge_order=1; dsamp=None means unit spacing."""
    n = tab.shape[1]
    f = tab.reshape(*[int(d) for d in dims], n)
    out = []
    for ax in range(len(dims)):
        h = 1.0 if dsamp is None else float(dsamp[ax])
        g = torch.empty_like(f)
        fm = torch.movedim(f, ax, 0)
        gm = torch.movedim(g, ax, 0)
        if fm.shape[0] < 2:
            raise ValueError("Shape of array too small to calculate a numerical gradient, "
                             "at least (edge_order + 1) elements are required.")
        gm[1:-1] = div(fm[2:] - fm[:-2], 2.0 * h)
        gm[0] = div(fm[1] - fm[0], h)
        gm[-1] = div(fm[-1] - fm[-2], h)
        out.append(g)
    return out


def amplitudes(ndims, ds, dr):
    """Geometrical spreading with the 1e-2 max-distance guard (K:316-330)."""
    maxdist = 1e-2 * (ds.max() + dr.max())
    if ndims == 2:
        return 1.0 / torch.sqrt(ds + maxdist), 1.0 / torch.sqrt(dr + maxdist), float(maxdist)
    return 1.0 / (ds + maxdist), 1.0 / (dr + maxdist), float(maxdist)


def angles(ndims, gs, gr, ni):
    """Angles with the vertical (K:364-393).  In 3D the receiver angle takes
    the sign of the source-side gradient, broadcast like the reference (so
    ns must equal nr, or one of them be 1)."""
    ns, nr = gs[0].shape[-1], gr[0].shape[-1]
    if ndims == 2:
        return (torch.atan2(gs[0], gs[1]).reshape(ni, ns),
                torch.atan2(gr[0], gr[1]).reshape(ni, nr))
    gs3 = torch.stack(gs)
    gr3 = torch.stack(gr)
    sgn = torch.sign(gs3[1])
    angs = sgn * torch.acos(gs3[-1] / torch.sqrt((gs3 ** 2).sum(0)))
    try:
        angr = sgn * torch.acos(gr3[-1] / torch.sqrt((gr3 ** 2).sum(0)))
        angr = angr.reshape(ni, nr)
    except RuntimeError as e:
        raise ValueError(f"operands could not be broadcast together ({e})") from None
    return angs.reshape(ni, ns), angr


def source_receiver_indices(ndims, x, srcs, recs):
    """six/rix: grid x-index of each trace's source and receiver (K:289-301)."""
    ns, nr = srcs.shape[1], recs.shape[1]
    if ndims == 2:
        dx = x[1] - x[0]
        six = np.tile((srcs[0] - x[0]) // dx, (nr, 1)).T.astype(int).ravel()
        rix = np.tile((recs[0] - x[0]) // dx, (ns, 1)).astype(int).ravel()
        return six, rix
    return np.zeros(nr * ns), np.zeros(nr * ns)


def aperture_izmin(six, rix, nz, apmax):
    """Smallest depth index iz for which |six - rix| / (iz + 1) < apmax, per
    trace (nz if none).  The offset-aperture taper of K:867-880 always picks
    aperturetap[0] = 1, so the aperture reduces to this depth gate; the ratio
    is evaluated in float64 exactly as the reference does."""
    off = np.abs(np.asarray(six, dtype=np.float64) - np.asarray(rix, dtype=np.float64))
    ratio = off[:, None] / (np.arange(nz, dtype=np.float64)[None, :] + 1)
    passing = ratio < apmax
    iz = np.where(passing.any(axis=1), passing.argmax(axis=1), nz)
    return iz.astype(np.int32)


def side_weights(amp, ang, angmin, angmax, aperturetap, scale=None):
    """Per-side factors of the dynamic weight
        2 cos((ts - tr)/2) As Ar / v * tap_s * tap_r
      = Ps Pr + Qs Qr,   P = A cos(t/2) tap,  Q = A sin(t/2) tap
    with 2/v folded into the source side (``scale``).  A side whose angle is
    not below angmax gets zero weight, which removes the triplet exactly like
    the reference's angle test (K:836-840).  The angle taper index is
    floor(20 (|t| - angmin) / dang) with Python's float floor division."""
    a = ang.abs()
    inside = a < angmax
    dang = angmax - angmin
    tap = torch.as_tensor(np.asarray(aperturetap, dtype=np.float64), device=ang.device)
    if dang > 0:
        q = div(20 * (a - angmin), dang, rounding_mode="floor")
        idx = torch.nan_to_num(q, nan=0.0).clamp(0, TAPER_LEN - 1).long()
        taper = torch.where((a >= angmin) & inside, tap[idx], torch.ones_like(a))
    else:
        taper = torch.ones_like(a)
    half = 0.5 * ang
    P = amp * torch.cos(half) * taper
    Q = amp * torch.sin(half) * taper
    if scale is not None:
        P = P * scale[:, None]
        Q = Q * scale[:, None]
    zero = torch.zeros((), dtype=P.dtype, device=P.device)
    return torch.where(inside, P, zero), torch.where(inside, Q, zero)


# ---------------------------------------------------------------------------
# device layout
# ---------------------------------------------------------------------------
def default_tile(ndims, nz):
    """Tile shape (tz, tx, ty) with tz*tx*ty = 64."""
    if ndims == 2:
        return (16, 4, 1)
    return (16, 2, 2)


def make_geom_plan(nx, ny, nz, tile):
    tz, tx, ty = tile
    if tz * tx * ty != 64:
        raise ValueError(f"tile {tile} must hold 64 points")
    p = _lib.KxPlan()
    p.abi_version = _lib.KX_ABI_VERSION
    p.nx, p.ny, p.nz = nx, ny, nz
    p.tz, p.tx, p.ty = tz, tx, ty
    p.ntz = -(-nz // tz)
    p.ntx = -(-nx // tx)
    p.nty = -(-ny // ty)
    p.ntiles = p.ntz * p.ntx * p.nty
    return p


def encode(plan, tab, scale, fmt, lo=-math.inf, hi=math.inf):
    """tab: float64 device tensor (ni, n) in reference order -> (off, base)
    device tensors in the tiled layout, offsets in samples (tab*scale).

    Entries outside [lo, hi) samples are ones that no trace sample can take
    (the caller derives the bounds from the other side's extreme times);
    they are passed to the encoder as NaN, which gives them an offset far
    outside every trace and leaves them out of their tile's base, so that
    unreachable points (huge times) cost the others no precision."""
    n = tab.shape[1]
    if not bool(torch.isfinite(tab).all()):
        raise ValueError("traveltime table contains non-finite values")
    # scale (1/dt) > 0: compare the table itself, no scaled copy
    out = (tab < lo / scale) | (tab >= hi / scale)
    if bool(out.any()):
        tab = torch.where(out, torch.full_like(tab, math.nan), tab)
    tabT = tab.t().contiguous()
    tdt = {_lib.KX_F32: torch.float32, _lib.KX_F16: torch.float16,
           _lib.KX_F64: torch.float64}[fmt]
    off = torch.empty((plan.ntiles, n, 64), dtype=tdt, device=tab.device)
    base = torch.empty((plan.ntiles, n), dtype=torch.int32, device=tab.device)
    rc = _lib.load().kx_encode_table(ctypes.byref(plan), _lib.ptr(tabT), n, float(scale), fmt,
                                     _lib.ptr(off), _lib.ptr(base), _lib.stream_ptr(tab.device))
    _lib.check(rc, "kx_encode_table")
    return off, base


def tile(plan, tab, fmt):
    """float64 (ni, n) -> [ntiles][n][64] of dtype fmt, zero padded."""
    n = tab.shape[1]
    tabT = tab.t().contiguous()
    tdt = {_lib.KX_F64: torch.float64, _lib.KX_F16: torch.float16}.get(fmt, torch.float32)
    out = torch.empty((plan.ntiles, n, 64), dtype=tdt, device=tab.device)
    rc = _lib.load().kx_tile_table(ctypes.byref(plan), _lib.ptr(tabT), n, fmt, _lib.ptr(out),
                                   _lib.stream_ptr(tab.device))
    _lib.check(rc, "kx_tile_table")
    return out


def weight_typical(wp, wq):
    """mean over the table's nonzero entries of |(wp, wq)|, in float64
    (0 when every entry is zero)."""
    wp, wq = wp.reshape(-1), wq.reshape(-1)
    total = torch.zeros((), dtype=torch.float64, device=wp.device)
    count = torch.zeros((), dtype=torch.float64, device=wp.device)
    step = 1 << 26
    for i in range(0, wp.numel(), step):
        a, b = wp[i:i + step].double(), wq[i:i + step].double()
        r = torch.sqrt(a * a + b * b)
        total += r.sum()
        count += (r > 0).sum()
    return float(total / count) if float(count) > 0 else 0.0


def weight_bound(wp, wq):
    """max over the table of |(wp, wq)|, in float64 (device tensors of one
    side's weight tables in the tiled layout)."""
    wp, wq = wp.reshape(-1), wq.reshape(-1)
    best = torch.zeros((), dtype=torch.float64, device=wp.device)
    step = 1 << 26
    for i in range(0, wp.numel(), step):
        a, b = wp[i:i + step].double(), wq[i:i + step].double()
        best = torch.maximum(best, torch.sqrt(a * a + b * b).max())
    return float(best)

