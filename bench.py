"""Kirchhoff forward + adjoint throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload 3d|2d|2dsmall|batch|autograd]
                    [--no-cpu] [--no-solve] [--no-extra] [--calib-copy]

A step is one forward (spread + wavelet convolution) of a dense image and
one adjoint (wavelet correlation + gather) of the observed data, i.e. the
operator pair one CGLS/LSQR iteration applies.  The dense image is the
migrated image K^T d (what the solver's forward sees after its first
adjoint), so no tile of the forward is skipped as empty.

Workloads (synthetic inputs, fixed seeds):
  3d       3D, byot straight-ray traveltimes for v(z) = 1800 + 0.6 z m/s;
           grid 120 x 128 x 128 at 10 m; nt=1001 at 4 ms; 5x5 sources and
           20x20 receivers on the surface; kinematic.          (default)
           The line's value is the float32-table step; the float16
           offset-table step, CGLS (20 iterations) s/solve for both, their
           errors against the float64 operator and the CPU baseline sit
           beside it, and under "extra" the 2d line and the batch.
  2d       2D, analytic constant velocity 2500 m/s, dynamic=True; grid
           nz=300 x nx=601 at 5 m; nt=1501 at 2 ms; 61 sources and 301
           receivers on the surface; 25 Hz Ricker; 8 dipping reflectors and
           point diffractors; LSQR and CGLS 30 iterations.
  2dsmall  2D, 2000 m/s, kinematic; 100 x 101 at 10 m; nt=651 at 4 ms;
           11 sources, 51 receivers.
  batch    64 independent 2D problems of uneven sizes (nz U[64,256], nx
           U[64,512], nsrc U[4,32], nrec U[32,256], nt U[256,1024] at 4 ms,
           velocity U[1500,3500] m/s, 15-30 Hz Ricker), fp16 tables; a step
           is one batched forward + one batched adjoint of all problems;
           also 20 CGLS iterations per problem, batched and one by one.
  autograd the 2d operator as a torch autograd Function on 16 models: loss
           0.5|Km-d|^2 + 1e-3 TV(m), 5% noise, 50 Adam steps (a step is
           the forward of all 16 models + backward + Adam, --adam-steps), and a
           central-difference check of the gradient along 32 directions.

Output: one JSON line.  value = 2 * ni * ns * nr / (t_fwd + t_adj) / 1e9
(Gsamples/s, one sample = one image point x source x receiver triplet).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Kirchhoff fwd+adj Gsamples/s, CGLS s/solve; FP16-table vs FP32 at ≤1e-3 err"
HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4  # MI355X peak engine clock


def ricker(t, f0):
    t = np.asarray(t)
    if len(t) % 2 == 0:
        t = t[:-1]
    tt = np.concatenate((np.flipud(-t[1:]), t))
    a = (np.pi * f0 * tt) ** 2
    w = (1 - 2 * a) * np.exp(-a)
    return w, int(np.argmax(np.abs(w)))


def workload_2d(rng):
    nx, nz, nt = 601, 300, 1501
    dx = dz = 5.0
    x = np.arange(nx) * dx
    z = np.arange(nz) * dz
    t = np.arange(nt) * 0.002
    srcs = np.vstack((np.linspace(x[0], x[-1], 61), np.zeros(61)))
    recs = np.vstack((np.linspace(x[0], x[-1], 301), np.zeros(301)))
    wav, wc = ricker(t[:41], 25.0)
    refl = np.zeros((nx, nz))
    X, Z = np.meshgrid(x, z, indexing="ij")
    for _ in range(8):
        dip = np.deg2rad(rng.uniform(-20, 20))
        z0 = rng.uniform(0.1, 0.9) * z[-1]
        zr = z0 + np.tan(dip) * (X - x[nx // 2])
        refl[np.abs(Z - zr) < dz / 2] += rng.uniform(-1, 1)
    for _ in range(20):
        refl[rng.integers(0, nx), rng.integers(nz // 5, nz)] += rng.uniform(-1, 1)
    kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=2500.0, wav=wav.astype(np.float32),
              wavcenter=wc, mode="analytic", dynamic=True)
    desc = dict(workload="2d-dynamic-601x300-61x301-nt1501", nx=nx, nz=nz, ns=61, nr=301,
                nt=nt, dynamic=True)
    return kw, refl, desc


def straight_ray_vz(pts_xyz, src_xyz, v0=1800.0, k=0.6):
    """Straight-ray traveltime through v(z) = v0 + k z: t = L ln(v2/v1)/(k dz),
    or L / v when dz = 0."""
    d = pts_xyz - src_xyz
    L = np.sqrt((d ** 2).sum(-1))
    z1 = src_xyz[..., 2]
    z2 = pts_xyz[..., 2]
    v1 = v0 + k * z1
    v2 = v0 + k * z2
    dzz = z2 - z1
    with np.errstate(divide="ignore", invalid="ignore"):
        tt = np.where(np.abs(dzz) > 1e-9, L * np.log(v2 / v1) / (k * dzz), L / v1)
    return tt


def workload_3d(rng, device):
    import torch
    ny = nx = 128
    nz = 120
    d = 10.0
    y = np.arange(ny) * d
    x = np.arange(nx) * d
    z = np.arange(nz) * d
    t = np.arange(1001) * 0.004
    sy, sx = np.meshgrid(np.linspace(y[0], y[-1], 5), np.linspace(x[0], x[-1], 5), indexing="ij")
    ry, rx = np.meshgrid(np.linspace(y[0], y[-1], 20), np.linspace(x[0], x[-1], 20),
                         indexing="ij")
    srcs = np.vstack((sy.ravel(), sx.ravel(), np.zeros(25)))
    recs = np.vstack((ry.ravel(), rx.ravel(), np.zeros(400)))
    wav, wc = ricker(t[:41], 20.0)
    # tables in float64 on the device: (ni, n), reference order (y, x, z)
    Y, X, Z = torch.meshgrid(*(torch.as_tensor(a, dtype=torch.float64, device=device)
                               for a in (y, x, z)), indexing="ij")
    P = torch.stack((Y.reshape(-1), X.reshape(-1), Z.reshape(-1)), -1)

    def table(pts):
        S = torch.as_tensor(pts.T, dtype=torch.float64, device=device)  # (n, 3) as (y, x, z)
        diff = P[:, None, :] - S[None, :, :]
        L = torch.sqrt((diff ** 2).sum(-1))
        v1 = 1800.0 + 0.6 * S[None, :, 2]
        v2 = 1800.0 + 0.6 * P[:, None, 2]
        dzz = P[:, None, 2] - S[None, :, 2]
        safe = torch.where(dzz.abs() > 1e-9, dzz, torch.ones_like(dzz))
        return torch.where(dzz.abs() > 1e-9, L * torch.log(v2 / v1) / (0.6 * safe), L / v1)

    ts = table(srcs)
    tr = table(recs)
    refl = np.zeros((ny, nx, nz))
    for _ in range(6):
        iz = rng.integers(10, nz - 5)
        refl[:, :, iz] = rng.uniform(-1, 1)
    refl += 0.01 * rng.standard_normal(refl.shape)
    kw = dict(z=z, x=x, y=y, t=t, srcs=srcs, recs=recs, vel=1800.0,
              wav=wav.astype(np.float32), wavcenter=wc, mode="byot", trav=(ts, tr))
    desc = dict(workload="3d-byot-vz-128x128x120-25x400-nt1001", nx=nx, ny=ny, nz=nz, ns=25,
                nr=400, nt=1001, dynamic=False)
    return kw, refl.reshape(ny * nx, nz), desc


def workload_2dsmall(rng):
    x = np.arange(101) * 10.0
    z = np.arange(100) * 10.0
    t = np.arange(651) * 0.004
    srcs = np.vstack((np.linspace(x[0], x[-1], 11), np.zeros(11)))
    recs = np.vstack((np.linspace(x[0], x[-1], 51), np.zeros(51)))
    wav, wc = ricker(t[:41], 20.0)
    refl = 0.01 * rng.standard_normal((101, 100))
    for iz in rng.integers(10, 90, 3):
        refl[:, iz] += rng.uniform(-1, 1)
    kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=2000.0, wav=wav.astype(np.float32),
              wavcenter=wc, mode="analytic", dynamic=False)
    desc = dict(workload="2d-kinematic-101x100-11x51-nt651", nx=101, nz=100, ns=11, nr=51,
                nt=651, dynamic=False)
    return kw, refl, desc


def measured_traffic(workload, kernel, tables="f32"):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summaries in profiles/ (a
    'TRAFFIC <workload> <kernel> <bytes> [tables]' line), with the file it
    came from, or (None, None)."""
    import glob
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*.txt"))):
        with open(path) as f:
            for line in f:
                m = re.match(r"TRAFFIC (\S+) (\S+) (\d+)(?: (\S+))?", line)
                if (m and m.group(1) == workload and m.group(2) == kernel
                        and m.group(4) in (None, tables)):
                    return int(m.group(3)), os.path.join("profiles", os.path.basename(path))
    return None, None


def algorithmic_bytes(desc, ni, table_bytes, dtype_bytes=4):
    """Separable-stream model: the receiver table streamed once per source,
    source tables once, the image once and the data once; dynamic tables
    (traveltime, two weights) triple the table terms, plus the velocity.
    Float16 tables (table_bytes 2) keep the source offsets in float32."""
    ns, nr, nt = desc["ns"], desc["nr"], desc["nt"]
    nk = 3 if desc["dynamic"] else 1
    sb = 4 if table_bytes == 2 else table_bytes
    b = nk * (ns * nr * ni * table_bytes + ns * ni * sb)
    b += ni * dtype_bytes + ns * nr * nt * dtype_bytes
    if desc["dynamic"]:
        b += ni * dtype_bytes
    return b


def compulsory_bytes(Op, dtype_bytes=4):
    """Bytes one kernel launch (spread or gather) must move at least: every
    table element, weight and base once (tile-padded device layout), the
    image once and the traces once; dynamic plans add the two weight tables
    per side and the per-trace aperture gate."""
    plan = Op.plan
    ntiles = int(plan.ntiles)
    ns, nr = int(plan.ns), int(plan.nr)
    tb = 8 if Op.dtype == np.float64 else 4
    tr = 2 if Op.tabledtype == np.float16 else tb  # float16: receiver offsets only
    b = (ns * tb + nr * tr) * ntiles * 64 + (ns + nr) * ntiles * 4
    if Op.dynamic:
        # two weight tables per side; float16 plans keep the receiver's in
        # float16
        b += 2 * (ns * dtype_bytes + nr * (2 if tr == 2 else dtype_bytes)) * ntiles * 64
        b += Op.nsnr * 4
    return b + Op.ni * dtype_bytes + Op.nsnr * Op.nt * dtype_bytes


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(kw, desc, m, d, target_s=2.0, runs=3):
    """The oracle (C/OpenMP restatement of the reference kernels, float32
    operator arithmetic, the reference's loop nests: forward parallel over
    sources, adjoint over image points) timed on a bounded sample of the
    workload -- the full image and receiver set with a subset of the
    sources -- at all host cores and at one thread: one warm-up run, then
    the median of `runs` timed forward + adjoint runs."""
    from oracle import oracle as O
    import torch
    L = O.lib()
    kwo = {k: v for k, v in kw.items()}
    if kwo.get("mode") == "byot":
        kwo["trav"] = tuple(a.cpu().numpy() if isinstance(a, torch.Tensor) else a
                            for a in kwo["trav"])
    kwo["dtype"] = "float32"
    nsrc_total = kwo["srcs"].shape[1]
    nall = int(L.or_num_threads())
    dd_all = np.asarray(d).reshape(nsrc_total, -1)
    mm = np.asarray(m, dtype=np.float32).ravel()

    def make(nsrc):
        k2 = dict(kwo)
        k2["srcs"] = kwo["srcs"][:, :nsrc]
        if k2.get("mode") == "byot":
            k2["trav"] = (np.ascontiguousarray(kwo["trav"][0][:, :nsrc]), kwo["trav"][1])
        return O.KirchhoffOracle(**k2), dd_all[:nsrc].ravel()

    def once(orc, dd):
        t0 = time.perf_counter()
        orc.matvec(mm)
        orc.rmatvec(dd)
        return time.perf_counter() - t0

    res = {}
    try:
        for label, nthr in (("all_cores", nall), ("one_thread", 1)):
            L.or_set_num_threads(nthr)
            # sources of the sample: a multiple of the thread count (the
            # forward runs one source per thread), sized from a probe run to
            # about target_s per run
            n1 = min(nthr, nsrc_total)
            orc, dd = make(n1)
            t1 = once(orc, dd)
            nsrc = n1
            if t1 < target_s / 2:
                nsrc = int(min(nsrc_total, n1 * target_s / max(t1, 1e-3)))
                nsrc = max(n1, (nsrc // n1) * n1)
                if nsrc != n1:
                    orc, dd = make(nsrc)
            once(orc, dd)  # warm-up
            ts = [once(orc, dd) for _ in range(runs)]
            t = float(np.median(ts))
            trip = 2.0 * orc.ni * nsrc * orc.nr
            res[label] = dict(value=round(trip / t / 1e9, 4), threads=nthr, sources=nsrc,
                              s_runs=[round(x, 3) for x in ts])
    finally:
        L.or_set_num_threads(nall)
    a = res["all_cores"]
    return dict(value=a["value"], unit="Gsamples/s", cores=nall, kind="port",
                cpu=cpu_model(), one_thread=res["one_thread"], all_cores=a,
                sample=f"{desc['workload']}: all {orc.nr} receivers and the full image "
                       f"with the first {a['sources']} of {nsrc_total} sources at "
                       f"{nall} threads ({res['one_thread']['sources']} at one "
                       f"thread); one forward + one adjoint (float32 operator "
                       f"arithmetic, float64 tables), median of {runs} after a warm-up")


def build_workload(name, dev):
    rng = np.random.default_rng(1003 if name == "3d" else 1002)
    if name == "2d":
        return workload_2d(rng)
    if name == "3d":
        return workload_3d(rng, dev)
    return workload_2dsmall(rng)


def time_operator(kw, refl, tables, dev, steps, warmup):
    """A step = kx_forward of the dense image K^T d (the spread with the
    wavelet convolution in the same launch, plus its two-block scale
    prepass) + the wavelet correlation of d + kx_gather, i.e. the operator
    pair of one solver iteration, each part bracketed by HIP events on the
    launch stream.  Returns the timings, the operator and the observed
    data."""
    import ctypes
    import torch
    import pylops_amd
    from pylops_amd import _lib
    tdt = "float16" if tables == "f16" else None
    Op = pylops_amd.Kirchhoff(dtype="float32", tabledtype=tdt, **kw)
    plan = Op.plan
    L = _lib.load()
    pp = ctypes.byref(plan)
    st = _lib.stream_ptr(dev)
    mt = torch.as_tensor(np.asarray(refl, dtype=np.float32).ravel(), device=dev)
    nd = Op.nsnr * Op.nt
    dobs = torch.empty(nd, dtype=torch.float32, device=dev)
    work = torch.empty(nd, dtype=torch.float32, device=dev)
    dcorr = torch.empty(nd, dtype=torch.float32, device=dev)
    img = torch.empty(Op.ni, dtype=torch.float32, device=dev)
    _lib.check(L.kx_forward(pp, _lib.ptr(mt), _lib.ptr(dobs), _lib.ptr(work), st))
    _lib.check(L.kx_adjoint(pp, _lib.ptr(dobs), _lib.ptr(img), _lib.ptr(work), st))
    mdense = img.clone()
    dout = torch.empty_like(dobs)
    h, hs = Op._dev_tables["h"], Op._dev_tables["hstar"]

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        _lib.check(L.kx_forward(pp, _lib.ptr(mdense), _lib.ptr(dout), _lib.ptr(work), st))
        if ev is not None:
            ev[1].record()
        _lib.check(L.kx_conv(_lib.ptr(dobs), _lib.ptr(dcorr), Op.nsnr, Op.nt, _lib.ptr(hs),
                             hs.numel(), _lib.KX_F32, st))
        if ev is not None:
            ev[2].record()
        _lib.check(L.kx_gather(pp, _lib.ptr(dcorr), _lib.ptr(img), st))
        if ev is not None:
            ev[3].record()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_spread = float(np.mean([e[0].elapsed_time(e[1]) for e in evs])) * 1e-3
    t_conv = float(np.mean([e[1].elapsed_time(e[2]) for e in evs])) * 1e-3
    t_gather = float(np.mean([e[2].elapsed_time(e[3]) for e in evs])) * 1e-3

    trip = float(Op.ni) * Op.ns * Op.nr
    res = {"tables": tables, "ms_per_step": round(wall / steps * 1e3, 4),
           "gsamples_s": round(2.0 * trip / (wall / steps) / 1e9, 3),
           "data_gsamples_s": round(2.0 * nd / (wall / steps) / 1e9, 4),
           # spread: kx_forward (the spread launch with the wavelet
           # convolution inside it, and its scale prepass)
           "kernel_ms": {"spread": round(t_spread * 1e3, 4),
                         "correlation": round(t_conv * 1e3, 4),
                         "gather": round(t_gather * 1e3, 4)},
           "launch_shapes": {"rblock": int(plan.rblock), "sorder": int(plan.sorder),
                             "sside": int(plan.sside), "gsplit": int(plan.gsplit),
                             "gtiles": int(plan.gtiles), "tile": list(Op.tile)},
           "compulsory_bytes": int(compulsory_bytes(Op))}
    return res, Op, dobs


def roofline(res, desc, ni, traffic_key):
    """Roofline of the dominant kernel of a step: achieved = compulsory
    bytes (every table element, base and weight once, image and traces
    once) over the kernel's mean HIP-event time, against the 8 TB/s HBM
    peak; the separable-stream model (receiver table once per source) and
    the rocprofv3 counter traffic beside it."""
    km = res["kernel_ms"]
    dom = "gather" if km["gather"] >= km["spread"] else "spread"
    t = km[dom] * 1e-3
    comp = res["compulsory_bytes"]
    sep = algorithmic_bytes(desc, ni, 2 if res["tables"] == "f16" else 4)
    traffic, src = measured_traffic(traffic_key, dom, res["tables"])
    out = {"bound": "hbm", "kernel": dom, "achieved": round(comp / t / 1e9, 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(comp / t / 1e9 / HBM_PEAK_GBS, 4),
           "traffic": traffic, "model": "compulsory", "model_bytes": int(comp),
           "separable_stream_bytes": int(sep),
           "separable_stream_frac": round(sep / t / 1e9 / HBM_PEAK_GBS, 4)}
    if traffic is not None:
        out["traffic_frac"] = round(traffic / t / 1e9 / HBM_PEAK_GBS, 4)
        out["traffic_source"] = src
    if dom == "spread":
        # what binds the spread: its two 64-bit LDS atomics per (source,
        # receiver, point) triplet, as lanes per clock per CU at the
        # published peak engine clock
        import torch
        cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
        atomics = 2.0 * desc["ns"] * desc["nr"] * ni
        out["lds_atomics"] = {"per_launch": atomics, "cus": cus, "clock_ghz": CLOCK_GHZ,
                              "lanes_per_clk_per_cu": round(atomics / t / (cus * CLOCK_GHZ * 1e9),
                                                            3)}
    return out


def operator_line(dev, args, name, cpu=True, precision=True, solve=True):
    """The bench line of one operator workload: float32 tables (the
    headline value) and float16 offset tables beside them, solver times,
    errors against the float64 operator and the CPU baseline."""
    import torch
    kw, refl, desc = build_workload(name, dev)
    r32, Op, dobs = time_operator(kw, refl, "f32", dev, args.steps, args.warmup)
    out = {
        "metric": METRIC,
        "value": r32["gsamples_s"],
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {**desc, "tables": "f32",
                   "step": "forward (spread + wavelet) of K^T d + adjoint (correlation + "
                           "gather) of d",
                   "kernel_ms": r32["kernel_ms"], "launch_shapes": r32["launch_shapes"],
                   "data_gsamples_s": r32["data_gsamples_s"]},
        "roofline": roofline(r32, desc, Op.ni, name),
    }
    niter = 20 if name == "3d" else (30 if name == "2d" else 10)
    if solve and not args.no_solve:
        out["solve"] = {"f32": time_solvers(Op, dobs, niter=niter,
                                            which=("cgls",) if name == "3d" else ("lsqr", "cgls"))}
    cpu_line = None
    if cpu and not args.no_cpu:
        cpu_line = cpu_baseline(kw, desc, refl, dobs.cpu().numpy())
    del Op, dobs
    torch.cuda.empty_cache()
    if name != "2dsmall":
        r16, Op16, dobs16 = time_operator(kw, refl, "f16", dev, args.steps, args.warmup)
        out["config"]["fp16_tables"] = {
            **r16, "speedup_vs_fp32": round(r32["ms_per_step"] / r16["ms_per_step"], 3),
            "spread_speedup": round(r32["kernel_ms"]["spread"] / r16["kernel_ms"]["spread"], 3),
            "gather_speedup": round(r32["kernel_ms"]["gather"] / r16["kernel_ms"]["gather"], 3),
            "roofline": roofline(r16, desc, Op16.ni, name)}
        if solve and not args.no_solve:
            out["solve"]["f16"] = time_solvers(
                Op16, dobs16, niter=niter, which=("cgls",) if name == "3d" else ("lsqr", "cgls"))
        del Op16, dobs16
        torch.cuda.empty_cache()
    if precision and not args.no_solve and name != "2dsmall":
        out["precision"] = precision_report(dev, name, niter=niter)
    if cpu_line is not None:
        out["cpu_baseline"] = cpu_line
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="3d", choices=["3d", "2d", "2dsmall", "batch", "autograd"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the solver timings and the precision reports (both run solvers)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads' lines (2D dynamic, 64-problem batch) that "
                         "the 3d line carries under 'extra'")
    ap.add_argument("--adam-steps", type=int, default=50,
                    help="Adam steps of the autograd workload")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 4 GiB device copy first (reference kernel for "
                         "FETCH_SIZE / WRITE_SIZE counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    import torch
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calib_copy:
        src = torch.ones(1024 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "batch":
        print(json.dumps(batch_line(dev, args)), flush=True)
        return
    if args.workload == "autograd":
        print(json.dumps(autograd_line(dev, args, nsteps=args.adam_steps)), flush=True)
        return
    out = operator_line(dev, args, args.workload)
    if args.workload == "3d" and not args.no_extra:
        out["extra"] = {"2d": operator_line(dev, args, "2d", cpu=False)}
        if not args.no_solve:
            out["extra"]["batch"] = time_batch(dev)[0]
    print(json.dumps(out), flush=True)


def time_solvers(Op, dobs, niter=30, runs=3, which=("cgls", "lsqr")):
    """Wall time of CGLS and / or LSQR solves of `niter` iterations on the
    observed data, inputs resident on the device, iterations replayed from
    a HIP graph (tol = atol = btol = 0, so every iteration runs): the median
    of `runs` solves, each run listed (the first timed CGLS solve of a
    process is at times ~40 ms slower).  The warm-up solve runs past the
    replay loop's first flag check."""
    import torch
    from pylops_amd.optimization.basic import cgls, lsqr
    res = {"niter": niter}
    for name, fn, kw in (("cgls", cgls, dict(tol=0.0)),
                         ("lsqr", lsqr, dict(atol=0.0, btol=0.0, calc_var=False))):
        if name not in which:
            continue
        fn(Op, dobs, niter=10, graph=True, **kw)  # past the replay loop's first flag check
        torch.cuda.synchronize()
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            r = fn(Op, dobs, niter=niter, graph=True, **kw)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        res[f"{name}_s_per_solve"] = round(float(np.median(times)), 4)
        res[f"{name}_s_runs"] = [round(t, 4) for t in times]
        res[f"{name}_iters"] = int(r[2])
    return res


def workload_batch(rng, nprob=64):
    """64 independent 2D target-oriented LSM problems of uneven sizes: nz
    U[64,256], nx U[64,512] at 10 m, nsrc U[4,32], nrec U[32,256] on the
    surface, nt U[256,1024] at 4 ms, constant velocity U[1500,3500] m/s,
    15-30 Hz Ricker; reflectivity = random layers plus 1% sparse spikes."""
    probs = []
    for _ in range(nprob):
        nz = int(rng.integers(64, 257))
        nx = int(rng.integers(64, 513))
        ns = int(rng.integers(4, 33))
        nr = int(rng.integers(32, 257))
        nt = int(rng.integers(256, 1025))
        x = np.arange(nx) * 10.0
        z = np.arange(nz) * 10.0
        t = np.arange(nt) * 0.004
        srcs = np.vstack((np.linspace(x[0], x[-1], ns), np.zeros(ns)))
        recs = np.vstack((np.linspace(x[0], x[-1], nr), np.zeros(nr)))
        wav, wc = ricker(t[:41], float(rng.uniform(15.0, 30.0)))
        kw = dict(z=z, x=x, t=t, srcs=srcs, recs=recs, vel=float(rng.uniform(1500.0, 3500.0)),
                  wav=wav.astype(np.float32), wavcenter=wc, mode="analytic", dynamic=False)
        refl = np.zeros((nx, nz), dtype=np.float32)
        nl = int(rng.integers(3, 9))
        for iz, a in zip(np.sort(rng.choice(np.arange(8, nz - 4), nl, replace=False)),
                         rng.uniform(-1, 1, nl)):
            refl[:, iz] = a
        spikes = rng.random((nx, nz)) < 0.01
        refl[spikes] += rng.uniform(-1, 1, int(spikes.sum())).astype(np.float32)
        probs.append((kw, refl))
    return probs


def batch_model_bytes(ops):
    """Separable-stream bytes of one application of every problem of a
    batch with fp16 offset tables (2 B per entry plus 4 B bases per tile)."""
    b = 0
    for op in ops:
        desc = {"ns": op.ns, "nr": op.nr, "nt": op.nt, "dynamic": op.dynamic}
        b += algorithmic_bytes(desc, op.ni, 2)
        ntiles = -(-op.ni // 64)
        b += (op.ns + op.nr) * ntiles * 4
    return b


def time_batch(dev, steps=5, warmup=2, niter=20, sequential=True):
    """The batch of 64 uneven problems with fp16 tables: batched forward +
    adjoint (persistent launches, HIP events around each), and 20 CGLS
    iterations per problem, all problems at once (graph-replayed
    iterations) and, for comparison, one problem after the other with the
    single-problem solver."""
    import torch
    import pylops_amd
    from pylops_amd.optimization.basic import cgls, cgls_batch
    from pylops_amd.waveeqprocessing import KirchhoffBatch, autotune_batch
    rng = np.random.default_rng(1004)
    probs = workload_batch(rng)
    kws = [dict(kw, dtype="float32", tabledtype="float16") for kw, _ in probs]
    # launch shapes from the autotuner (measured on this device, not cached)
    t0 = time.perf_counter()
    tuned = autotune_batch(kws, reps=2, cache=False)
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    ops = [pylops_amd.Kirchhoff(**kw, tile=tuned["tile"]) for kw in kws]
    B = KirchhoffBatch(ops, rblock=tuned["rblock"], sorder=tuned["sorder"],
                       gsplit=tuned["gsplit"], gtiles=tuned["gtiles"])
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    m = torch.as_tensor(np.concatenate([r.ravel() for _, r in probs]), device=dev)
    y = torch.empty(B.shape[0], dtype=torch.float32, device=dev)
    a = torch.empty(B.shape[1], dtype=torch.float32, device=dev)
    dout = torch.empty_like(y)
    B._into(m, y, False)
    B._into(y, a, True)   # dense model: the migrated image
    for _ in range(warmup):
        B._into(a, dout, False)
        B._into(y, a, True)
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for e in evs:
        e[0].record()
        B._into(a, dout, False)
        e[1].record()
        B._into(y, a, True)
        e[2].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    t_fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in evs])) * 1e-3
    t_adj = float(np.mean([e[1].elapsed_time(e[2]) for e in evs])) * 1e-3
    trip = float(sum(op.ni * op.ns * op.nr for op in ops))
    res = {"workload": "lsm_batch_64_2d_uneven", "nprob": len(ops), "tables": "f16",
           "rblock": B.rblock, "gsplit": B.gsplit, "gtiles": B.gtiles, "sorder": B.sorder,
           "spread_items": int(B._batch.n_spread), "gather_items": int(B._batch.n_gather),
           "model_points": int(B.shape[1]), "data_samples": int(B.shape[0]),
           "triplets": trip, "setup_s": round(t_setup, 3),
           "autotune": {"tile": list(tuned["tile"]), "rblock": tuned["rblock"],
                        "sorder": tuned["sorder"], "gsplit": tuned["gsplit"],
                        "gtiles": tuned["gtiles"],
                        "ms_fwd_adj": tuned["ms"], "tuning_s": round(t_tune, 2)},
           "ms_per_step": round(wall * 1e3, 4),
           "fwd_ms": round(t_fwd * 1e3, 4), "adj_ms": round(t_adj * 1e3, 4),
           "fwd_adj_gsamples_s": round(2 * trip / wall / 1e9, 2),
           "model_bytes": int(batch_model_bytes(ops)),
           "compulsory_bytes": int(sum(compulsory_bytes(op) for op in ops))}
    B._into(m, y, False)
    cgls_batch(B, y, niter=2, tol=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = cgls_batch(B, y, niter=niter, tol=0.0)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    res.update({"niter": niter, "cgls_s_per_batch_solve": round(t_batch, 4),
                "iters": sorted(set(int(r[2]) for r in sol))})
    if sequential:
        ys = B.split_data(y)
        t0 = time.perf_counter()
        for op, yp in zip(ops, ys):
            cgls(op, yp, niter=niter, tol=0.0, graph=True)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
        # the single-problem adjoint picks its own tile groups and split of
        # a group's sources over waves; with the batch's it sums in the same
        # order, and the solutions are bitwise those of the batch
        same = True
        for p in range(4):
            ops[p].gsplit, ops[p].gtiles = B.gsplit, B.gtiles
            xs = cgls(ops[p], ys[p], niter=niter, tol=0.0, graph=True)[0]
            same &= bool(torch.equal(xs.reshape(-1), sol[p][0].reshape(-1)))
            ops[p].gsplit = ops[p].gtiles = 0
        res.update({"cgls_s_sequential_solves": round(t_seq, 4),
                    "speedup_vs_sequential": round(t_seq / t_batch, 3),
                    "bitwise_equal_to_sequential_same_split": same})
    return res, probs


def cpu_baseline_batch(probs, budget_s=12.0):
    """The oracle on a bounded sample of the batch: problems in order, one
    forward + one adjoint each (float32 operator arithmetic, fp64 tables),
    until the time budget is spent."""
    from oracle import oracle as O
    nthreads = O.lib().or_num_threads()
    rng = np.random.default_rng(0)
    trip, tsum, k = 0.0, 0.0, 0
    for kw, refl in probs:
        kwo = dict(kw)
        kwo["dtype"] = "float32"
        orc = O.KirchhoffOracle(**kwo)
        d = rng.standard_normal(orc.ns * orc.nr * orc.nt).astype(np.float32)
        t0 = time.perf_counter()
        orc.matvec(refl.ravel())
        orc.rmatvec(d)
        tsum += time.perf_counter() - t0
        trip += 2.0 * orc.ni * orc.ns * orc.nr
        k += 1
        if tsum > budget_s:
            break
    return dict(value=trip / tsum / 1e9, unit="Gsamples/s", cores=int(nthreads), kind="port",
                sample=f"lsm_batch_64_2d_uneven, the first {k} of {len(probs)} problems, "
                       f"one forward + one adjoint each (float32 operator arithmetic), "
                       f"{tsum:.1f} s")


def batch_line(dev, args):
    res, probs = time_batch(dev, steps=args.steps, warmup=args.warmup,
                            sequential=not args.no_solve)
    t_dom = max(res["fwd_ms"], res["adj_ms"]) * 1e-3
    dom = "batch_forward" if res["fwd_ms"] >= res["adj_ms"] else "batch_adjoint"
    comp = res["compulsory_bytes"]
    achieved = comp / t_dom / 1e9
    traffic, src = measured_traffic(res["workload"], dom, "f16")
    out = {
        "metric": METRIC,
        "value": res["fwd_adj_gsamples_s"],
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": res["workload"], "tables": "f16",
                   "step": "batched forward of K^T d + batched adjoint of d, 64 problems",
                   "batch": res},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "model": "compulsory", "model_bytes": comp,
                     "separable_stream_bytes": res["model_bytes"],
                     "separable_stream_frac": round(res["model_bytes"] / t_dom / 1e9
                                                    / HBM_PEAK_GBS, 4)},
    }
    if traffic is not None:
        out["roofline"]["traffic_frac"] = round(traffic / t_dom / 1e9 / HBM_PEAK_GBS, 4)
        out["roofline"]["traffic_source"] = src
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_batch(probs)
    return out


def isotropic_tv(m, eps=1e-8):
    """sum of sqrt(dx^2 + dz^2 + eps) over a batch of (nx, nz) models,
    forward differences (torch autograd)."""
    dx = m[:, 1:, :-1] - m[:, :-1, :-1]
    dz = m[:, :-1, 1:] - m[:, :-1, :-1]
    return (dx * dx + dz * dz + eps).sqrt().sum()


def autograd_line(dev, args, nbatch=16, nsteps=50, ndir=32):
    """The 2D dynamic operator (601x300, 61x301, nt=1501) as a torch
    autograd Function on a batch of 16 models: loss = 0.5 |(K m - d)/s|^2 +
    1e-3 TV(m) against data from 16 seeded ground-truth models with 5%
    Gaussian noise (s = the clean data's standard deviation), 50 Adam steps
    from zero.  The gradient of the data term (the adjoint kernel) is
    checked against central differences along 32 random unit directions
    with step 1e-3 |m|.  One step = forward of all 16 models + adjoint of
    all 16 residuals (the operator's matmat / rmatmat: one launch of each
    kernel per model) + TV and Adam."""
    import torch
    import pylops_amd
    from pylops_amd import TorchOperator
    rng = np.random.default_rng(1005)
    kw, _, desc = workload_2d(rng)
    refls = [workload_2d(rng)[1] for _ in range(nbatch)]
    Op = pylops_amd.Kirchhoff(dtype="float32", **kw)
    Top = TorchOperator(Op, batch=True, device="gpu")
    nx, nz = Op.nx, Op.nz
    mtrue = torch.as_tensor(np.stack([r.reshape(nx * nz) for r in refls]), device=dev,
                            dtype=torch.float32)
    with torch.no_grad():
        dclean = Top.apply(mtrue)
        sig = float(dclean.std())
        g = torch.Generator(device="cpu").manual_seed(5)
        noise = torch.randn(dclean.shape, generator=g).to(dev)
        dobs = (dclean + 0.05 * sig * noise) / sig
    lam = 1e-3

    def data_term(m):
        r = Top.apply(m) / sig - dobs
        return 0.5 * torch.sum(r * r, dtype=torch.float64)

    def loss_fn(m):
        return data_term(m) + lam * isotropic_tv(m.view(-1, nx, nz)).double()

    m = torch.zeros_like(mtrue, requires_grad=True)
    opt = torch.optim.Adam([m], lr=0.02)
    for _ in range(args.warmup):  # the batched operator and torch's kernels
        opt.zero_grad()
        loss_fn(m).backward()
    m.data.zero_()
    opt = torch.optim.Adam([m], lr=0.02)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(nsteps)]
    losses = []
    t0 = time.perf_counter()
    for k in range(nsteps):
        opt.zero_grad()
        ev[k][0].record()
        pred = Top.apply(m)
        ev[k][1].record()
        r = pred / sig - dobs
        # squares in float32, summed in float64 (no float64 copy of the
        # 16 x 27.6M residual samples)
        loss = (0.5 * torch.sum(r * r, dtype=torch.float64)
                + lam * isotropic_tv(m.view(-1, nx, nz)).double())
        loss.backward()
        ev[k][2].record()
        opt.step()
        losses.append(loss.detach())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_fwd = float(np.mean([e[0].elapsed_time(e[1]) for e in ev])) * 1e-3
    t_bwd = float(np.mean([e[1].elapsed_time(e[2]) for e in ev])) * 1e-3
    losses = [float(v) for v in losses]
    err_true = float((m.detach() - mtrue).norm() / mtrue.norm())
    # adjoint-kernel gradient vs central differences of the data term; the
    # difference of the two losses is summed as (r+ - r-)(r+ + r-) / 2 so that
    # it does not cancel
    m0 = m.detach().clone().requires_grad_(True)
    data_term(m0).backward()
    gr = m0.grad.double()
    h = 1e-3 * float(m0.detach().norm())
    errs = []
    with torch.no_grad():
        for _ in range(ndir):
            u = torch.as_tensor(rng.standard_normal(m0.shape), dtype=torch.float64, device=dev)
            u /= u.norm()
            rp = (Top.apply((m0.double() + h * u).float()) / sig - dobs).double()
            rm = (Top.apply((m0.double() - h * u).float()) / sig - dobs).double()
            fd = float((0.5 * (rp - rm) * (rp + rm)).sum() / (2 * h))
            an = float((gr * u).sum())
            errs.append(abs(fd - an) / max(abs(an), 1e-30))
    trip = float(Op.ni) * Op.ns * Op.nr * nbatch
    step_s = wall / nsteps
    bytes_alg = algorithmic_bytes(desc, Op.ni, 4) * nbatch
    out = {
        "metric": METRIC,
        "value": round(2.0 * trip / step_s / 1e9, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {**desc, "workload": "autograd_2d_dynamic_batch16_tv_adam",
                   "batch": nbatch, "tv_weight": lam, "noise": 0.05,
                   "optimizer": "Adam lr=0.02",
                   "step": "forward of 16 models + backward (adjoint of 16 residuals, TV) + Adam",
                   "forward_ms": round(t_fwd * 1e3, 3),
                   "loss_backward_ms": round(t_bwd * 1e3, 3),
                   "s_total": round(wall, 3),
                   "launch_shapes": {"rblock": Op.plan.rblock, "sorder": Op.plan.sorder,
                                     "gsplit": Op.plan.gsplit},
                   "loss_first": losses[0], "loss_last": losses[-1],
                   "model_rel_err_last": err_true,
                   "fd_check": {"step": "1e-3 |m|", "directions": ndir, "term": "data",
                                "median_rel_err": float(np.median(errs)),
                                "max_rel_err": float(np.max(errs))}},
        "roofline": {"bound": "hbm", "kernel": "forward of 16 models",
                     "achieved": round(bytes_alg / t_fwd / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(bytes_alg / t_fwd / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "model_bytes": int(bytes_alg)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(kw, desc, refls[0], dclean[0].cpu().numpy())
    return out


def precision_report(dev, which, niter):
    """Errors of the float32 operator with float32 and float16 tables
    against the float64 operator (float64 tables, float64 arithmetic; its
    kernels are checked against the oracle to 1e-10), for the forward of the
    reflectivity, the adjoint of the float64 data and the model after
    `niter` CGLS iterations; and the wall time of forward(reflectivity) +
    adjoint(data) with both table formats.  The 2D reflectivity is sparse
    and the spread skips waves whose image points are all zero, so there
    these times are shorter than the dense-image step of the headline line
    (the 3D reflectivity carries noise everywhere and is dense)."""
    import torch
    import pylops_amd
    from pylops_amd.optimization.basic import cgls
    rng = np.random.default_rng(1003 if which == "3d" else 1002)
    kw, refl, desc = workload_3d(rng, dev) if which == "3d" else workload_2d(rng)
    Op64 = pylops_amd.Kirchhoff(dtype="float64", **kw)
    m64 = torch.as_tensor(np.asarray(refl, dtype=np.float64).ravel(), device=dev)
    d64 = Op64 @ m64
    a64 = Op64.H @ d64
    r64 = cgls(Op64, d64, niter=niter, tol=0.0, graph=True)
    x64, cost64 = r64[0].reshape(-1), np.asarray(r64[5])
    del Op64
    torch.cuda.empty_cache()

    def rel(u, v):
        return float((u.double().reshape(-1) - v.reshape(-1)).norm() / v.norm())

    res = {"workload": desc["workload"], "reference": "float64 operator", "cgls_niter": niter}
    for tabs in ("f32", "f16"):
        Op = pylops_amd.Kirchhoff(dtype="float32", tabledtype="float16" if tabs == "f16" else None,
                                  **kw)
        m = m64.float()
        d32 = d64.float()
        d = Op @ m
        a = Op.H @ d32
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 3
        for _ in range(n):
            Op @ m
            Op.H @ d32
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / n
        rx = cgls(Op, d32, niter=niter, tol=0.0, graph=True)
        x, cost = rx[0].reshape(-1), np.asarray(rx[5])
        res[tabs] = {"ms_fwd_adj_refl": round(dt * 1e3, 3),
                     "rel_err_fwd": rel(d, d64), "rel_err_adj": rel(a, a64),
                     "rel_err_cgls_x": rel(x, x64),
                     "cgls_residual_norm_rel_diff_max": float(
                         np.max(np.abs(cost - cost64) / cost64)),
                     "cgls_residual_norm_reduction": float(cost[-1] / cost[0]),
                     "cgls_residual_norm_reduction_fp64": float(cost64[-1] / cost64[0])}
        del Op
        torch.cuda.empty_cache()
    res["fp16_speedup_refl"] = round(res["f32"]["ms_fwd_adj_refl"]
                                     / res["f16"]["ms_fwd_adj_refl"], 3)
    return res


if __name__ == "__main__":
    main()
