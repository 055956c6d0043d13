"""Measured choice of the Kirchhoff kernels' launch shapes.

The forward and adjoint kernels take four shape parameters that do not
change their results beyond rounding (the float32 forward not at all):

- ``tile`` (tz, tx, ty), the image block of one wavefront (the tables are
  rebuilt for each candidate);
- ``rblock``, receivers (or sources) per forward workgroup, ``sorder``,
  the forward's lane order, and ``sside``, which side a forward workgroup
  blocks (single problems);
- ``gsplit``, adjoint waves per tile.

:func:`autotune` times every candidate on the device with a dense random
model (every tile busy, as in solver iterations after the first) and
returns the fastest; :func:`autotune_batch` does the same for a batch,
where the tile applies to every problem and the other three to the batch's
persistent launches.  The forward and adjoint are tuned independently for
each tile (their shapes do not interact); the tile with the smallest sum
wins.  Results are kept in a JSON cache keyed by geometry, precision and
device name.
"""
import ctypes
import json
import os
import warnings

import numpy as np
import torch

from pylops_amd import _lib

__all__ = ["autotune", "autotune_batch", "TILES_2D", "TILES_3D"]

TILES_2D = ((16, 4, 1), (8, 8, 1), (32, 2, 1), (4, 16, 1))
TILES_3D = ((16, 2, 2), (8, 4, 2), (4, 4, 4), (8, 8, 1), (16, 4, 1))
RBLOCKS = (16, 8, 4, 2, 1)
SORDERS = (_lib.KX_ORDER_NATURAL, _lib.KX_ORDER_ROTATED)
SSIDES = (_lib.KX_SIDE_RECEIVERS, _lib.KX_SIDE_SOURCES)
GSPLITS = (1, 2, 4)
GTILES = (1, 2, 4)


def _cache_path(path):
    if path is not None:
        return path
    env = os.environ.get("PYLOPS_AMD_TUNE_CACHE")
    if env:
        return env
    return os.path.join(os.path.expanduser("~"), ".cache", "pylops_amd", "tune.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, key, value):
    """Adds one result to the cache file; a cache that cannot be written
    (read-only home, full disk) is skipped with a warning."""
    d = _load(path)
    d[key] = value
    try:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError as e:
        warnings.warn(f"autotune cache {path} not written: {e}")


def _geometry(kw):
    ndims = 2 if kw.get("y") is None else 3
    ny = 1 if ndims == 2 else len(kw["y"])
    ns = np.asarray(kw["srcs"]).shape[1]
    nr = np.asarray(kw["recs"]).shape[1]
    return ndims, len(kw["z"]), len(kw["x"]), ny, ns, nr, len(kw["t"])


def _key(kind, kws, device):
    dev = torch.cuda.get_device_name(device)
    parts = []
    for kw in kws:
        g = _geometry(kw)
        parts.append("x".join(str(v) for v in g) + f":{int(bool(kw.get('dynamic', False)))}"
                     f":{kw.get('dtype', 'float64')}:{kw.get('tabledtype')}")
    return f"{kind}|{dev}|" + ";".join(parts)


def _timer(fn, reps):
    """Average device time of fn() over reps runs after one warm-up."""
    fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def _candidates_fwd(dtype, dynamic, nt, sides=(_lib.KX_SIDE_RECEIVERS,), orders=SORDERS,
                    lds=65536):
    max_rb = 4 if dtype == "float64" else (8 if dynamic else 16)
    if dtype == "float64":
        sides = (_lib.KX_SIDE_RECEIVERS,)  # float64 operators block receivers
    for rb in RBLOCKS:
        # the forward's LDS rows: nt rounded up to 2 mod 16 samples of 8 bytes
        if rb > max_rb or (rb > 1 and rb * (nt + (2 - nt) % 16) * 8 > lds):
            continue  # more registers than two workgroups per CU, or LDS
        for so in orders if rb > 1 else (_lib.KX_ORDER_NATURAL,):
            for side in sides:
                yield rb, so, side


def autotune(reps=3, tiles=None, cache=None, **kw):
    """Fastest launch shapes for ``Kirchhoff(**kw)``: returns a dict with
    ``tile``, ``rblock``, ``sorder``, ``sside``, ``gsplit``, ``gtiles``
    (pass it on to ``Kirchhoff``) and the measured ``ms`` (forward kernel +
    adjoint kernel).  ``tiles`` overrides the candidate tiles; ``cache`` is the
    JSON cache file (default ``$PYLOPS_AMD_TUNE_CACHE`` or
    ``~/.cache/pylops_amd/tune.json``; ``False`` disables it)."""
    from pylops_amd.waveeqprocessing.kirchhoff import Kirchhoff
    _lib.require_device()
    kw = {k: v for k, v in kw.items()
          if k not in ("tile", "rblock", "sorder", "gsplit", "sside", "gtiles")}
    device = kw.get("device") or torch.device("cuda", torch.cuda.current_device())
    path = None if cache is False else _cache_path(cache)
    key = _key("op", [kw], device)
    if path is not None:
        hit = _load(path).get(key)
        if hit is not None:
            hit["tile"] = tuple(hit["tile"])
            hit.setdefault("sside", _lib.KX_SIDE_AUTO)  # entries from before sside was tuned
            hit.setdefault("gtiles", 0)  # entries from before gtiles was tuned
            return hit
    ndims = _geometry(kw)[0]
    tiles = tiles or (TILES_2D if ndims == 2 else TILES_3D)
    L = _lib.load()
    dtype = str(np.dtype(kw.get("dtype", "float64")))
    tdt = torch.float64 if dtype == "float64" else torch.float32
    best = None
    for tile in tiles:
        if tile[0] > len(kw["z"]) * 2:
            continue  # a tile twice as tall as the image wastes lanes
        Op = Kirchhoff(**kw, tile=tile)
        plan = Op.plan
        g = torch.Generator(device="cpu").manual_seed(0)
        m = torch.randn(Op.ni, generator=g, dtype=tdt).to(Op.device)
        traces = torch.empty(Op.nsnr * Op.nt, dtype=tdt, device=Op.device)
        img = torch.empty(Op.ni, dtype=tdt, device=Op.device)
        st = _lib.stream_ptr(Op.device)
        L.kx_spread(ctypes.byref(plan), _lib.ptr(m), _lib.ptr(traces), st)

        def spread():
            _lib.check(L.kx_spread(ctypes.byref(plan), _lib.ptr(m), _lib.ptr(traces), st),
                       "kx_spread")

        def gather():
            _lib.check(L.kx_gather(ctypes.byref(plan), _lib.ptr(traces), _lib.ptr(img), st),
                       "kx_gather")

        fwd = []
        lds = torch.cuda.get_device_properties(Op.device).shared_memory_per_block
        for rb, so, side in _candidates_fwd(dtype, kw.get("dynamic", False), len(kw["t"]),
                                            SSIDES, lds=lds):
            plan.rblock, plan.sorder, plan.sside = rb, so, side
            fwd.append((_timer(spread, reps), rb, so, side))
        adj = []
        for gs in GSPLITS:
            for gt in GTILES:
                plan.gsplit, plan.gtiles = gs, gt
                adj.append((_timer(gather, reps), gs, gt))
        tf, rb, so, side = min(fwd)
        ta, gs, gt = min(adj)
        cand = {"tile": tuple(tile), "rblock": rb, "sorder": so, "sside": side, "gsplit": gs,
                "gtiles": gt, "ms": round(tf + ta, 4), "ms_forward": round(tf, 4),
                "ms_adjoint": round(ta, 4)}
        if best is None or cand["ms"] < best["ms"]:
            best = cand
        del Op, m, traces, img
    if best is None:
        raise ValueError("no candidate tile fits this image")
    if path is not None:
        _store(path, key, best)
    return best


def autotune_batch(kws, reps=3, tiles=None, cache=None):
    """Fastest launch shapes for ``KirchhoffBatch([Kirchhoff(**kw) for kw
    in kws])``: returns ``tile`` (for every problem), ``rblock``,
    ``sorder``, ``gsplit``, ``gtiles`` (for the batch) and the measured
    ``ms`` (batched forward + batched adjoint)."""
    from pylops_amd.waveeqprocessing.kirchhoff import Kirchhoff
    from pylops_amd.waveeqprocessing.kirchhoffbatch import KirchhoffBatch
    _lib.require_device()
    kws = [{k: v for k, v in kw.items()
            if k not in ("tile", "rblock", "sorder", "gsplit", "gtiles")} for kw in kws]
    device = kws[0].get("device") or torch.device("cuda", torch.cuda.current_device())
    path = None if cache is False else _cache_path(cache)
    key = _key("batch", kws, device)
    if path is not None:
        hit = _load(path).get(key)
        if hit is not None:
            hit["tile"] = tuple(hit["tile"])
            hit.setdefault("gtiles", 0)  # entries from before gtiles was tuned
            return hit
    ndims = _geometry(kws[0])[0]
    tiles = tiles or (TILES_2D if ndims == 2 else TILES_3D)
    dtype = str(np.dtype(kws[0].get("dtype", "float64")))
    tdt = torch.float64 if dtype == "float64" else torch.float32
    dynamic = bool(kws[0].get("dynamic", False))
    best = None
    for tile in tiles:
        ops = [Kirchhoff(**kw, tile=tile) for kw in kws]
        B0 = KirchhoffBatch(ops)
        g = torch.Generator(device="cpu").manual_seed(0)
        m = torch.randn(B0.shape[1], generator=g, dtype=tdt).to(B0.device)
        d = torch.empty(B0.shape[0], dtype=tdt, device=B0.device)
        a = torch.empty(B0.shape[1], dtype=tdt, device=B0.device)
        B0._into(m, d, False)
        fwd = []
        for rb, so, _ in _candidates_fwd(dtype, dynamic, max(len(kw["t"]) for kw in kws),
                                         orders=(_lib.KX_ORDER_NATURAL,
                                                 _lib.KX_ORDER_ROTATED)):
            B = KirchhoffBatch(ops, rblock=rb, sorder=so)
            if B.rblock != rb:
                continue  # does not fit the batch's longest trace
            fwd.append((_timer(lambda: B._into(m, d, False), reps), rb, so))
        adj = []
        for gs in GSPLITS:
            for gt in (1, 2):
                B = KirchhoffBatch(ops, gsplit=gs, gtiles=gt)
                adj.append((_timer(lambda: B._into(d, a, True), reps), gs, gt))
        tf, rb, so = min(fwd)
        ta, gs, gt = min(adj)
        cand = {"tile": tuple(tile), "rblock": rb, "sorder": so, "gsplit": gs, "gtiles": gt,
                "ms": round(tf + ta, 4), "ms_forward": round(tf, 4), "ms_adjoint": round(ta, 4)}
        if best is None or cand["ms"] < best["ms"]:
            best = cand
        del ops, B0, m, d, a
    if path is not None:
        _store(path, key, best)
    return best
